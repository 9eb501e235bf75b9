// Semi-global matching over uint32 (SSD) and float (NCC) cost volumes,
// standard geometry: the optimiser half of the reference's SSDDev +
// SemiGlobalDev and CrossCorrelationDev + SemiGlobalDev pairings
// (semiglobaldev.cpp:16-28 takes any DSIMem; commonstereomatcherfactory.cpp:
// 39-81 composes any cost with it).
//
// Recurrence per direction (semiglobalcpu.cpp:87-125 with constant
// penalties), in this operation order:
//   t = min(L'(d), min(L'(d-1), L'(d+1)) + P1)
//   E = min(t - m, P2)            m = min_d L'(d)
//   L(p, d) = C(p, d) + E
// and L = C at a path start, or when every cell of L' is invalid (the path
// restarts there; the reference's float arithmetic would give inf - inf).
// Invalid operands drop out of the minima. uint32: invalid 0xFFFFFFFF and
// every addition saturates at it. float: +inf invalid (NaN and -inf costs
// read as +inf), IEEE round-to-nearest additions. The aggregated volume is
// S = L_first + L_second + ... over the enabled directions in increasing
// index, one direction per launch, so float sums have a fixed order.
//
// Mapping: one wave per line; lane l holds d = l*K + k (K = ceil(D/64) <= 4)
// in registers; the d-1 / d+1 neighbours across lanes move with whole-wave
// DPP shifts (the vacated end lane reads the invalid value); m is a DPP
// minimum over the wave. Costs stream through a prefetch ring of PF steps.
//
// uint32 volumes with P2 <= 255 (byte planes): since E = L - C lies in
// [0, P2] (0 at a path start or restart), every line of every enabled
// direction runs in one launch and stores E as one byte per cell in its
// direction's plane; a second kernel forms S = n*C + sum of the n planes'
// E, saturating at the invalid value (a sum of non-negative terms
// saturates to the same value in any order, and L = C + E saturates only
// when S does). That moves 5 bytes per cell and direction through the
// scans instead of 12 and fills the GPU with every direction's lines at
// once. With P2 <= 127 the scans read relative cost bytes instead of the
// uint32 costs: per pixel min(C - Cmin, 2 P2), 2 P2 on invalid cells, 0xFF
// on every cell of a pixel with no valid cell (read as invalid, restarting
// the path as the true costs do). Shifting a pixel's costs by a constant
// and clipping them at 2 P2 above its minimum changes no increment E (a
// cell at or above Cmin + 2 P2 stays at or above m + P2 with either cost),
// so the planes are the same; the sum still adds n times the true C. The
// scans then read 1 byte per cell and direction instead of 4.
// Float volumes whose planes stay below 2 GiB run the same way with
// a float plane of L per direction (the per-direction kernel's operations,
// cell for cell) and a second kernel that adds the planes in increasing
// direction index, which gives S the roundings of one launch per direction
// adding into it. Other volumes run one launch per direction adding into
// S.
#include <cstdlib>
#include <string>
#include <type_traits>

#include "sgm_device.hpp"

namespace tdvh {

template <typename T> struct WideOps;
template <> struct WideOps<uint32_t> {
    static constexpr int INV_BITS = -1;
    __device__ static uint32_t inv() { return INV32; }
    __device__ static uint32_t read(uint32_t v) { return v; }
    __device__ static uint32_t add(uint32_t a, uint32_t b) { return __builtin_elementwise_add_sat(a, b); }
    __device__ static uint32_t mn(uint32_t a, uint32_t b) { return min(a, b); }
    __device__ static bool valid(uint32_t v) { return v != INV32; }
    // min(t - m, P2) for t >= m (t invalid: t - m is above any penalty)
    __device__ static uint32_t excess(uint32_t t, uint32_t m, uint32_t p2) { return min(t - m, p2); }
};
template <> struct WideOps<float> {
    static constexpr int INV_BITS = 0x7F800000;
    __device__ static float inv() { return __builtin_inff(); }
    __device__ static float read(float v)
    {
        return (v < __builtin_inff() && v > -__builtin_inff()) ? v : __builtin_inff();
    }
    __device__ static float add(float a, float b) { return a + b; }
    __device__ static float mn(float a, float b) { return __builtin_fminf(a, b); }
    __device__ static bool valid(float v) { return v < __builtin_inff(); }
    __device__ static float excess(float t, float m, float p2) { return __builtin_fminf(t - m, p2); }
};

// DPP move of a 32-bit value; lanes without a source read the invalid value.
template <int CTRL, typename T> __device__ __forceinline__ T wide_dpp(T v)
{
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(WideOps<T>::INV_BITS, __builtin_bit_cast(int, v),
                                                             CTRL, 0xF, 0xF, false));
}

// Minimum over the wave, returned to every lane (wave-uniform).
template <typename T> __device__ __forceinline__ T wave_min(T v)
{
    typedef WideOps<T> O;
    v = O::mn(v, wide_dpp<0x128>(v)); // row_ror:8
    v = O::mn(v, wide_dpp<0x124>(v)); // row_ror:4
    v = O::mn(v, wide_dpp<0x122>(v)); // row_ror:2
    v = O::mn(v, wide_dpp<0x121>(v)); // row_ror:1
    const T r0 = __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
    const T r1 = __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
    const T r2 = __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
    const T r3 = __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
    return O::mn(O::mn(r0, r1), O::mn(r2, r3));
}

template <typename T> struct WideArgs {
    const T *C;
    int64_t csx, csy; // cost strides (elements)
    T *S;
    int64_t ssx, ssy; // sum strides (elements)
    int W, H, D;
    T P1, P2;
};

// Lines of direction `dir`, one wave each (4 per workgroup). FIRST: the
// direction writes S; otherwise it adds to it.
template <typename T, int K, int PF, bool FIRST>
__global__ __launch_bounds__(256) void sgm_wide_dir_kernel(const WideArgs<T> a, int dir)
{
    typedef WideOps<T> O;
    const int line = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int l = threadIdx.x & 63;
    if (line >= dir_lines(dir, a.W, a.H)) return; // whole waves
    int sx, sy, len;
    line_geom(dir, line, a.W, a.H, sx, sy, len);
    const int dx = dir_dx(dir), dy = dir_dy(dir);
    const int klast = len - 1;
    const int d0 = l * K;

    // costs (and the running sum) of step k, clamped to the line
    struct Slot {
        T c[K], s[K];
    };
    auto fetch = [&](Slot &sl, int k) {
        const int kk = k < klast ? k : klast;
        const int x = sx + kk * dx, y = sy + kk * dy;
        const T *cp = a.C + x * a.csx + y * a.csy + d0;
        const T *sp = a.S + x * a.ssx + y * a.ssy + d0;
#pragma unroll
        for (int q = 0; q < K; q++) {
            const bool in = d0 + q < a.D;
            sl.c[q] = in ? O::read(cp[q]) : O::inv();
            if constexpr (!FIRST) sl.s[q] = in ? sp[q] : O::inv();
        }
    };
    Slot ring[PF];
#pragma unroll
    for (int u = 0; u < PF; u++) fetch(ring[u], u);

    T Lp[K];
#pragma unroll
    for (int q = 0; q < K; q++) Lp[q] = O::inv();
    T m = O::inv(); // every cell of L' invalid: the next step starts the path

    auto body = [&](Slot &sl, int k) {
        T c[K], s[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            c[q] = sl.c[q];
            if constexpr (!FIRST) s[q] = sl.s[q];
        }
        fetch(sl, k + PF);
        T Ln[K];
        if (!O::valid(m)) { // wave-uniform
#pragma unroll
            for (int q = 0; q < K; q++) Ln[q] = c[q];
        } else {
            const T left = wide_dpp<0x138>(Lp[K - 1]);  // wave_shr:1
            const T right = wide_dpp<0x130>(Lp[0]);     // wave_shl:1
#pragma unroll
            for (int q = 0; q < K; q++) {
                const T lm = q == 0 ? left : Lp[q - 1];
                const T lp = q == K - 1 ? right : Lp[q + 1];
                const T t = O::mn(Lp[q], O::add(O::mn(lm, lp), a.P1));
                Ln[q] = O::add(c[q], O::excess(t, m, a.P2));
            }
        }
        T mn = Ln[0];
#pragma unroll
        for (int q = 1; q < K; q++) mn = O::mn(mn, Ln[q]);
        m = wave_min(mn);
        const int x = sx + k * dx, y = sy + k * dy;
        T *sp = a.S + x * a.ssx + y * a.ssy + d0;
#pragma unroll
        for (int q = 0; q < K; q++) {
            Lp[q] = Ln[q];
            if (d0 + q < a.D) {
                if constexpr (FIRST) sp[q] = Ln[q];
                else sp[q] = O::add(s[q], Ln[q]);
            }
        }
    };

    int k0 = 0;
    for (; k0 + PF <= len; k0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) body(ring[u], k0 + u);
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (k0 + u < len) body(ring[u], k0 + u);
}

template <typename T, int K>
static void launch_wide_k(const WideArgs<T> &a, int dir, bool first, hipStream_t s)
{
    constexpr int PF = K <= 2 ? 8 : 4;
    const int blocks = (dir_lines(dir, a.W, a.H) + 3) / 4;
    if (first) sgm_wide_dir_kernel<T, K, PF, true><<<blocks, 256, 0, s>>>(a, dir);
    else sgm_wide_dir_kernel<T, K, PF, false><<<blocks, 256, 0, s>>>(a, dir);
}

// ---- uint32 volumes, byte planes ------------------------------------------------

// Minimum over the wave, returned to every lane: row shifts 1, 2, 4, 8 leave
// each row's minimum in its lane 15, row_bcast:15 / :31 fold the rows into
// lane 63.
template <int CTRL, int ROWS> __device__ __forceinline__ uint32_t min_dpp_wide(uint32_t v)
{
    return min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, CTRL, ROWS, 0xF, false));
}
__device__ __forceinline__ uint32_t wave_min_u32_all(uint32_t v)
{
    v = min_dpp_wide<0x111, 0xF>(v);
    v = min_dpp_wide<0x112, 0xF>(v);
    v = min_dpp_wide<0x114, 0xF>(v);
    v = min_dpp_wide<0x118, 0xF>(v);
    v = min_dpp_wide<0x142, 0xA>(v);
    v = min_dpp_wide<0x143, 0xC>(v);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

struct WideEArgs {
    const uint8_t *C;   // uint32 cost volume
    int64_t csx, csy;   // its strides in bytes
    const uint8_t *CR;  // relative cost bytes, DP bytes per pixel ([y][x][d]), or null
    int DP;             // D rounded up to 4
    uint8_t *planes;    // one W*H*D byte plane per launched direction, [y][x][d]
    size_t plane_bytes;
    int W, H, D;
    uint32_t P1, P2;
    int ndirs;
    int dir[8];
    int begin[9]; // first wave of each direction's lines (the last: all waves)
};

// One wave per line, the lines of every launched direction in one grid
// (longest directions first). Costs of step k are read through a buffer
// resource based at the step's pixel (ranged to its D cells, or to nothing
// past the line), so the lane offsets stay constant and the prefetch needs
// no clamp.
template <int K, int PF, bool REL>
__global__ __launch_bounds__(64) void sgm_wide_e_kernel(const WideEArgs a)
{
    const int gw = blockIdx.x;
    int s = 0;
    while (s < a.ndirs - 1 && gw >= a.begin[s + 1]) s++; // wave-uniform
    const int dir = a.dir[s], line = gw - a.begin[s];
    int sx, sy, len;
    line_geom(dir, line, a.W, a.H, sx, sy, len);
    const int dx = dir_dx(dir), dy = dir_dy(dir);
    const int l = threadIdx.x, d0 = l * K;
    // REL: K bytes per lane at byte d0 of the pixel's DP bytes (aligned for
    // K = 1, 2, 4; K = 3 reads bytes)
    const uint32_t lane_off = REL ? (uint32_t)d0 : (uint32_t)d0 * 4;
    const uint32_t cellb = REL ? (uint32_t)a.DP : (uint32_t)a.D * 4;
    const int64_t stepb = REL ? ((int64_t)dy * a.W + dx) * a.DP : dx * a.csx + dy * a.csy;
    const uint8_t *fp = REL ? a.CR + ((int64_t)sy * a.W + sx) * a.DP : a.C + sx * a.csx + sy * a.csy;
    int kf = 0;
    auto fetch = [&](uint32_t(&slot)[K]) {
        const buf_t r = buf_rsrc(fp, kf < len ? cellb : 0u);
        if constexpr (!REL) {
            buf_load<K, 0, 2>(r, lane_off, slot); // read once: non-temporal
        } else if constexpr (K == 4) {
            slot[0] = __builtin_amdgcn_raw_buffer_load_b32(r, lane_off, 0, 0);
        } else if constexpr (K == 2) {
            slot[0] = __builtin_amdgcn_raw_buffer_load_b16(r, lane_off, 0, 0);
        } else {
#pragma unroll
            for (int q = 0; q < K; q++) slot[q] = __builtin_amdgcn_raw_buffer_load_b8(r, lane_off + q, 0, 0);
        }
        fp += stepb;
        kf++;
    };
    uint32_t ring[PF][K];
#pragma unroll
    for (int u = 0; u < PF; u++) fetch(ring[u]);

    const buf_t pr = buf_rsrc(a.planes + (size_t)s * a.plane_bytes, (uint32_t)a.plane_bytes);
    uint32_t eo = (uint32_t)(((size_t)sy * a.W + sx) * a.D) + (uint32_t)d0;
    const uint32_t de = (uint32_t)((dy * a.W + dx) * a.D);
    const int nvalid = a.D - d0; // cells of this lane below D (may be <= 0)

    uint32_t Lp[K];
#pragma unroll
    for (int q = 0; q < K; q++) Lp[q] = INV32;
    uint32_t m = INV32; // every cell of L' invalid: the next step starts the path

    auto body = [&](uint32_t(&slot)[K]) {
        uint32_t c[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            if constexpr (REL) {
                const uint32_t b = (K == 2 || K == 4) ? (slot[0] >> (8 * q)) & 0xFFu : slot[q] & 0xFFu;
                c[q] = q < nvalid && b != 0xFFu ? b : INV32;
            } else {
                c[q] = q < nvalid ? slot[q] : INV32;
            }
        }
        fetch(slot);
        uint32_t E[K], Ln[K];
        if (m == INV32) { // wave-uniform: path start or restart
#pragma unroll
            for (int q = 0; q < K; q++) E[q] = 0;
        } else {
            const uint32_t left = wide_dpp<0x138>(Lp[K - 1]); // wave_shr:1
            const uint32_t right = wide_dpp<0x130>(Lp[0]);    // wave_shl:1
#pragma unroll
            for (int q = 0; q < K; q++) {
                const uint32_t lm = q == 0 ? left : Lp[q - 1];
                const uint32_t lp = q == K - 1 ? right : Lp[q + 1];
                const uint32_t nb = lm < lp ? lm : lp;
                const uint32_t up = __builtin_elementwise_add_sat(nb, a.P1);
                const uint32_t t = Lp[q] < up ? Lp[q] : up;
                E[q] = (t - m) < a.P2 ? (t - m) : a.P2;
            }
        }
        uint32_t mn = INV32, e = 0;
#pragma unroll
        for (int q = 0; q < K; q++) {
            Ln[q] = __builtin_elementwise_add_sat(c[q], E[q]);
            Lp[q] = Ln[q];
            mn = min(mn, Ln[q]);
            e |= E[q] << (8 * q);
        }
        m = wave_min_u32_all(mn);
        if (nvalid >= K) {
            if constexpr (K == 4) __builtin_amdgcn_raw_buffer_store_b32(e, pr, eo, 0, 0);
            else if constexpr (K == 2) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)e, pr, eo, 0, 0);
            else if constexpr (K == 1) __builtin_amdgcn_raw_buffer_store_b8((unsigned char)e, pr, eo, 0, 0);
            else {
                __builtin_amdgcn_raw_buffer_store_b16((unsigned short)e, pr, eo, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b8((unsigned char)(e >> 16), pr, eo + 2, 0, 0);
            }
        } else {
#pragma unroll
            for (int q = 0; q < K; q++)
                if (q < nvalid)
                    __builtin_amdgcn_raw_buffer_store_b8((unsigned char)(e >> (8 * q)), pr, eo + q, 0, 0);
        }
        eo += de;
    };
    int k0 = 0;
    for (; k0 + PF <= len; k0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) body(ring[u]);
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (k0 + u < len) body(ring[u]);
}

// S(x, y, d) = n*C + sum of the n planes' E, saturating at the invalid value
// (invalid where C is); one thread per pixel and 4 consecutive d. VEC: D % 4
// == 0 and 16-byte aligned cost and sum cells (one 16-byte load of C, one
// 4-byte load per plane, one 16-byte store of S).
__device__ __forceinline__ uint32_t wide_s(uint32_t c, uint32_t esum, int n)
{
    const uint64_t v = (uint64_t)c * (uint64_t)n + esum;
    return c == INV32 ? INV32 : (uint32_t)(v < (uint64_t)INV32 ? v : (uint64_t)INV32);
}

template <bool VEC>
__global__ __launch_bounds__(256) void sgm_wide_sum_kernel(const uint8_t *__restrict__ C, int64_t csx,
                                                           int64_t csy, const uint8_t *__restrict__ planes,
                                                           size_t plane_bytes, int n, uint8_t *__restrict__ S,
                                                           int64_t ssx, int64_t ssy, int W, int H, int D)
{
    const int G = (D + 3) / 4;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)W * H * G) return;
    const int g = (int)(idx % G);
    const int64_t pix = idx / G;
    const int x = (int)(pix % W), y = (int)(pix / W);
    const int d = 4 * g;
    const uint8_t *cp = C + x * csx + y * csy + (int64_t)d * 4;
    uint8_t *sp = S + x * ssx + y * ssy + (int64_t)d * 4;
    const uint8_t *ep = planes + pix * D + d;
    if constexpr (VEC) {
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 c = __builtin_nontemporal_load((const u32x4 *)cp);
        uint32_t e0 = 0, e1 = 0, e2 = 0, e3 = 0;
        for (int r = 0; r < n; r++) {
            const uint32_t w = __builtin_nontemporal_load((const uint32_t *)(ep + (size_t)r * plane_bytes));
            e0 += w & 0xFF;
            e1 += (w >> 8) & 0xFF;
            e2 += (w >> 16) & 0xFF;
            e3 += w >> 24;
        }
        const u32x4 out = {wide_s(c.x, e0, n), wide_s(c.y, e1, n), wide_s(c.z, e2, n), wide_s(c.w, e3, n)};
        *(u32x4 *)sp = out;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (d + q >= D) break;
            uint32_t esum = 0;
            for (int r = 0; r < n; r++) esum += ep[(size_t)r * plane_bytes + q];
            *(uint32_t *)(sp + 4 * q) = wide_s(*(const uint32_t *)(cp + 4 * q), esum, n);
        }
    }
}

// Order-preserving keys of the lowest-index WTA over a sum: uint32 as is
// (the invalid value all ones), float non-finite sums all ones.
__device__ __forceinline__ uint32_t wide_key(uint32_t v) { return v; }
__device__ __forceinline__ uint32_t wide_key(float v)
{
    if (!(v < __builtin_inff())) return 0xFFFFFFFFu;
    const uint32_t b = __builtin_bit_cast(uint32_t, v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// The lowest index holding the minimum key over a pixel's P lanes (each lane
// keeps its first minimum over increasing d; ties between lanes go to the
// lower d), or 0 when no cell is below the invalid key: wtacpu.cpp:20-33.
template <int P>
__device__ __forceinline__ int pixel_argmin(uint32_t bv, uint32_t bd)
{
#pragma unroll
    for (int off = P / 2; off > 0; off >>= 1) {
        const uint32_t ov = (uint32_t)__shfl_xor((int)bv, off, 64);
        const uint32_t od = (uint32_t)__shfl_xor((int)bd, off, 64);
        if (ov < bv || (ov == bv && od < bd)) {
            bv = ov;
            bd = od;
        }
    }
    return bv == 0xFFFFFFFFu ? 0 : (int)bd;
}

// Threads per pixel of the fused sum + WTA: the power of two at or above
// ceil(D / 4) (D <= 256).
static int sum_wta_lanes(int D)
{
    const int g = (D + 3) / 4;
    int p = 1;
    while (p < g) p *= 2;
    return p;
}

// Plane sum + lowest-index WTA without writing S (the caller asked only for
// disparities): P threads per pixel (the power of two at or above
// ceil(D / 4)), 4 cells each, the sums of sgm_wide_sum_kernel.
template <int P, bool VEC>
__global__ __launch_bounds__(256) void sgm_wide_sum_wta_kernel(const uint8_t *__restrict__ C, int64_t csx,
                                                               int64_t csy, const uint8_t *__restrict__ planes,
                                                               size_t plane_bytes, int n, int W, int H, int D,
                                                               int16_t *__restrict__ disp, int pitch)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t pix = idx / P;
    const int t = (int)(idx % P);
    const bool live = pix < (int64_t)W * H; // every lane joins the shuffles
    const int x = live ? (int)(pix % W) : 0, y = live ? (int)(pix / W) : 0;
    const int d = 4 * t;
    uint32_t bv = 0xFFFFFFFFu, bd = 0xFFFFFFFFu;
    if (live && d < D) {
        const uint8_t *cp = C + x * csx + y * csy + (int64_t)d * 4;
        const uint8_t *ep = planes + pix * D + d;
        uint32_t c[4], e[4] = {0, 0, 0, 0};
        if constexpr (VEC) {
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            const u32x4 cv = __builtin_nontemporal_load((const u32x4 *)cp);
            c[0] = cv.x, c[1] = cv.y, c[2] = cv.z, c[3] = cv.w;
            for (int r = 0; r < n; r++) {
                const uint32_t w = __builtin_nontemporal_load((const uint32_t *)(ep + (size_t)r * plane_bytes));
#pragma unroll
                for (int q = 0; q < 4; q++) e[q] += (w >> (8 * q)) & 0xFFu;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                c[q] = d + q < D ? *(const uint32_t *)(cp + 4 * q) : INV32;
                if (d + q < D)
                    for (int r = 0; r < n; r++) e[q] += ep[(size_t)r * plane_bytes + q];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t k = wide_key(wide_s(c[q], e[q], n));
            if (d + q < D && k < bv) {
                bv = k;
                bd = (uint32_t)(d + q);
            }
        }
    }
    const int best = pixel_argmin<P>(bv, bd);
    if (live && t == 0) disp[(size_t)y * pitch + x] = (int16_t)best;
}

// The same with 16 cells per thread (D % 16 == 0, 16-byte aligned cells):
// 16-byte loads of C and of every plane. N = 8: the eight planes' loads and
// C's are all issued before any sum (the plane count a compile-time
// constant); N = 0: a loop over the n planes of a direction mask.
template <int P, int N>
__global__ __launch_bounds__(256) void sgm_wide_sum_wta16_kernel(const uint8_t *__restrict__ C, int64_t csx,
                                                                 int64_t csy, const uint8_t *__restrict__ planes,
                                                                 size_t plane_bytes, int n, int W, int H, int D,
                                                                 int16_t *__restrict__ disp, int pitch)
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t pix = idx / P;
    const int t = (int)(idx % P);
    const bool live = pix < (int64_t)W * H; // every lane joins the shuffles
    const int x = live ? (int)(pix % W) : 0, y = live ? (int)(pix / W) : 0;
    const int d = 16 * t;
    uint32_t bv = 0xFFFFFFFFu, bd = 0xFFFFFFFFu;
    if (live && d < D) {
        const uint8_t *cp = C + x * csx + y * csy + (int64_t)d * 4;
        const uint8_t *ep = planes + pix * D + d;
        uint32_t e[16];
#pragma unroll
        for (int q = 0; q < 16; q++) e[q] = 0;
        u32x4 c4[4];
        if constexpr (N > 0) {
            u32x4 w[N];
#pragma unroll
            for (int r = 0; r < N; r++)
                w[r] = __builtin_nontemporal_load((const u32x4 *)(ep + (size_t)r * plane_bytes));
#pragma unroll
            for (int g = 0; g < 4; g++)
                c4[g] = __builtin_nontemporal_load((const u32x4 *)(cp + 16 * g));
#pragma unroll
            for (int r = 0; r < N; r++)
#pragma unroll
                for (int q = 0; q < 16; q++) e[q] += (w[r][q / 4] >> (8 * (q % 4))) & 0xFFu;
        } else {
            for (int r = 0; r < n; r++) {
                const u32x4 w = __builtin_nontemporal_load((const u32x4 *)(ep + (size_t)r * plane_bytes));
#pragma unroll
                for (int q = 0; q < 16; q++) e[q] += (w[q / 4] >> (8 * (q % 4))) & 0xFFu;
            }
#pragma unroll
            for (int g = 0; g < 4; g++)
                c4[g] = __builtin_nontemporal_load((const u32x4 *)(cp + 16 * g));
        }
        const int nn = N > 0 ? N : n;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const u32x4 c = c4[g];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t k = wide_key(wide_s(c[q], e[4 * g + q], nn));
                if (k < bv) {
                    bv = k;
                    bd = (uint32_t)(d + 4 * g + q);
                }
            }
        }
    }
    const int best = pixel_argmin<P>(bv, bd);
    if (live && t == 0) disp[(size_t)y * pitch + x] = (int16_t)best;
}

// Relative cost bytes of a uint32 volume (the form above): P threads per
// pixel (the power of two at or above ceil(D / 4)), 4 cells each (VEC: one
// 16-byte load, D % 4 == 0 and 16-byte aligned cells); the pixel's minimum
// by shuffles within its P lanes.
template <int P, bool VEC>
__global__ __launch_bounds__(256) void sgm_wide_rel_kernel(const uint8_t *__restrict__ C, int64_t csx,
                                                           int64_t csy, int W, int H, int D, int DP,
                                                           uint32_t cap, uint8_t *__restrict__ CR)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t pix = idx / P;
    const int t = (int)(idx % P);
    const bool live = pix < (int64_t)W * H; // every lane joins the shuffles
    const int x = live ? (int)(pix % W) : 0, y = live ? (int)(pix / W) : 0;
    const uint32_t *cp = (const uint32_t *)(C + x * csx + y * csy);
    uint32_t v[4];
    if (VEC && live && 4 * t < D) {
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 w4 = __builtin_nontemporal_load((const u32x4 *)(cp + 4 * t));
        v[0] = w4.x, v[1] = w4.y, v[2] = w4.z, v[3] = w4.w;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int d = 4 * t + q;
            v[q] = live && d < D ? cp[d] : INV32;
        }
    }
    uint32_t mn = min(min(v[0], v[1]), min(v[2], v[3]));
#pragma unroll
    for (int off = P / 2; off > 0; off >>= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, off, 64));
    uint32_t w = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint32_t b = mn == INV32 ? 0xFFu : (v[q] == INV32 ? cap : min(v[q] - mn, cap));
        w |= b << (8 * q);
    }
    if (live && 4 * t < DP) *(uint32_t *)(CR + pix * DP + 4 * t) = w;
}

template <int K>
static void launch_wide_e_k(const WideEArgs &a, hipStream_t s)
{
    constexpr int PF = 8;
    if (a.CR) sgm_wide_e_kernel<K, PF, true><<<a.begin[a.ndirs], 64, 0, s>>>(a);
    else sgm_wide_e_kernel<K, PF, false><<<a.begin[a.ndirs], 64, 0, s>>>(a);
}

// Whether the scans read relative cost bytes: P2 <= 127 (every byte but
// the marker below 255) and the bytes below 2 GiB; TDVH_SGM_WIDE_REL=0 reads
// the uint32 costs.
static bool wide_rel_wanted(int W, int H, int D, uint32_t p2)
{
    static const int mode = [] {
        const char *e = getenv("TDVH_SGM_WIDE_REL");
        return e ? atoi(e) : 1;
    }();
    return mode != 0 && p2 <= 127 && (size_t)W * H * ((D + 3) / 4 * 4) < ((size_t)1 << 31);
}

// The byte-plane form for uint32 volumes (P2 <= 255, planes below 2 GiB).
static int run_wide_e(tdvh_ctx *ctx, const tdvh_dsi *C, const tdvh_sgm_params *p, const tdvh_dsi *S,
                      int16_t *disp, int disp_pitch, hipStream_t s)
{
    const int W = C->width, H = C->height, D = C->depth;
    WideEArgs a{};
    a.C = (const uint8_t *)C->ptr;
    a.csx = C->stride_x * 4;
    a.csy = C->stride_y * 4;
    a.plane_bytes = (size_t)W * H * D;
    a.W = W;
    a.H = H;
    a.D = D;
    a.P1 = (uint32_t)p->p1;
    a.P2 = (uint32_t)p->p2;
    // longest lines first: horizontal, vertical, then the diagonals
    static const int order[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    const bool wide = W >= H;
    int ord[8];
    for (int i = 0; i < 8; i++) ord[i] = order[i];
    if (!wide) { ord[0] = 2; ord[1] = 3; ord[2] = 0; ord[3] = 1; }
    a.begin[0] = 0;
    for (int i = 0; i < 8; i++) {
        const int dir = ord[i];
        if (!(p->dir_mask & (1 << dir))) continue;
        a.dir[a.ndirs] = dir;
        a.begin[a.ndirs + 1] = a.begin[a.ndirs] + dir_lines(dir, W, H);
        a.ndirs++;
    }
    if (a.ndirs == 0) return TDVH_OK;
    void *pb = nullptr;
    TDVH_TRY(ctx->scratch("sgm_wide_planes", a.plane_bytes * a.ndirs, &pb));
    a.planes = (uint8_t *)pb;
    a.DP = (D + 3) / 4 * 4;
    if (wide_rel_wanted(W, H, D, a.P2)) {
        void *cr = nullptr;
        TDVH_TRY(ctx->scratch("sgm_wide_rel", (size_t)W * H * a.DP, &cr));
        a.CR = (const uint8_t *)cr;
        const int g = (D + 3) / 4;
        const int P = g <= 1 ? 1 : g <= 2 ? 2 : g <= 4 ? 4 : g <= 8 ? 8 : g <= 16 ? 16 : g <= 32 ? 32 : 64;
        const int64_t threads = (int64_t)W * H * P;
        const unsigned blocks = (unsigned)((threads + 255) / 256);
        const uint32_t cap = 2 * a.P2;
        prof_begin(s);
        const bool vec = D % 4 == 0 && a.csx % 16 == 0 && a.csy % 16 == 0 &&
                         ((uintptr_t)C->ptr & 15) == 0;
        switch (P) {
#define CASE(p)                                                                                    \
    case p:                                                                                        \
        if (vec) sgm_wide_rel_kernel<p, true><<<blocks, 256, 0, s>>>(a.C, a.csx, a.csy, W, H, D, a.DP, cap, (uint8_t *)cr); \
        else sgm_wide_rel_kernel<p, false><<<blocks, 256, 0, s>>>(a.C, a.csx, a.csy, W, H, D, a.DP, cap, (uint8_t *)cr); \
        break;
            CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
        }
        TDVH_TRY(check_launch("sgm_wide_rel_kernel", s));
    }
    prof_begin(s);
    switch ((D + 63) / 64) {
    case 1: launch_wide_e_k<1>(a, s); break;
    case 2: launch_wide_e_k<2>(a, s); break;
    case 3: launch_wide_e_k<3>(a, s); break;
    default: launch_wide_e_k<4>(a, s); break;
    }
    TDVH_TRY(check_launch("sgm_wide_lines_kernel", s));
    if (!S) { // disparities only: the sum feeds the WTA directly
        const bool cvec = D % 4 == 0 && C->stride_x % 4 == 0 && C->stride_y % 4 == 0 &&
                          ((uintptr_t)C->ptr & 15) == 0;
        if (cvec && D % 16 == 0) {
            const int P16 = sum_wta_lanes(D / 4); // lanes of 16 cells
            const unsigned b16 = (unsigned)(((int64_t)W * H * P16 + 255) / 256);
            prof_begin(s);
            switch (P16) {
#define CASE(q)                                                                                    \
    case q:                                                                                        \
        if (a.ndirs == 8) sgm_wide_sum_wta16_kernel<q, 8><<<b16, 256, 0, s>>>(a.C, a.csx, a.csy, a.planes, a.plane_bytes, a.ndirs, W, H, D, disp, disp_pitch); \
        else sgm_wide_sum_wta16_kernel<q, 0><<<b16, 256, 0, s>>>(a.C, a.csx, a.csy, a.planes, a.plane_bytes, a.ndirs, W, H, D, disp, disp_pitch); \
        break;
                CASE(1) CASE(2) CASE(4) CASE(8) CASE(16)
#undef CASE
            }
            return check_launch("sgm_wide_sum_wta_kernel", s);
        }
        const int P = sum_wta_lanes(D);
        const unsigned wblocks = (unsigned)(((int64_t)W * H * P + 255) / 256);
        prof_begin(s);
        switch (P) {
#define CASE(q)                                                                                    \
    case q:                                                                                        \
        if (cvec) sgm_wide_sum_wta_kernel<q, true><<<wblocks, 256, 0, s>>>(a.C, a.csx, a.csy, a.planes, a.plane_bytes, a.ndirs, W, H, D, disp, disp_pitch); \
        else sgm_wide_sum_wta_kernel<q, false><<<wblocks, 256, 0, s>>>(a.C, a.csx, a.csy, a.planes, a.plane_bytes, a.ndirs, W, H, D, disp, disp_pitch); \
        break;
            CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
        }
        return check_launch("sgm_wide_sum_wta_kernel", s);
    }
    const int64_t threads = (int64_t)W * H * ((D + 3) / 4);
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    const bool vec = D % 4 == 0 && C->stride_x % 4 == 0 && C->stride_y % 4 == 0 &&
                     S->stride_x % 4 == 0 && S->stride_y % 4 == 0 && ((uintptr_t)C->ptr & 15) == 0 &&
                     ((uintptr_t)S->ptr & 15) == 0;
    prof_begin(s);
    if (vec)
        sgm_wide_sum_kernel<true><<<blocks, 256, 0, s>>>(a.C, a.csx, a.csy, a.planes, a.plane_bytes,
                                                         a.ndirs, (uint8_t *)S->ptr, S->stride_x * 4,
                                                         S->stride_y * 4, W, H, D);
    else
        sgm_wide_sum_kernel<false><<<blocks, 256, 0, s>>>(a.C, a.csx, a.csy, a.planes, a.plane_bytes,
                                                          a.ndirs, (uint8_t *)S->ptr, S->stride_x * 4,
                                                          S->stride_y * 4, W, H, D);
    return check_launch("sgm_wide_sum_kernel", s);
}

// ---- float volumes, float path-cost planes -----------------------------------------

struct WideFArgs {
    const uint8_t *C;   // float cost volume
    int64_t csx, csy;   // its strides in bytes
    uint8_t *planes;    // one W*H*D float plane per launched direction, [y][x][d]
    size_t plane_bytes;
    int W, H, D;
    float P1, P2;
    int ndirs;
    int dir[8];
    int plane[8]; // plane of each launched direction: its rank by direction index
    int begin[9]; // first wave of each direction's lines (the last: all waves)
};

// One wave per line, the lines of every launched direction in one grid
// (longest directions first), as sgm_wide_e_kernel; each step's L, computed
// as in sgm_wide_dir_kernel, goes to the direction's plane.
template <int K, int PF>
__global__ __launch_bounds__(64) void sgm_wide_l_kernel(const WideFArgs a)
{
    typedef WideOps<float> O;
    const int gw = blockIdx.x;
    int s = 0;
    while (s < a.ndirs - 1 && gw >= a.begin[s + 1]) s++; // wave-uniform
    const int dir = a.dir[s], line = gw - a.begin[s];
    int sx, sy, len;
    line_geom(dir, line, a.W, a.H, sx, sy, len);
    const int dx = dir_dx(dir), dy = dir_dy(dir);
    const int l = threadIdx.x, d0 = l * K;
    const uint32_t lane_off = (uint32_t)d0 * 4, cellb = (uint32_t)a.D * 4;
    const int64_t stepb = dx * a.csx + dy * a.csy;
    const uint8_t *fp = a.C + sx * a.csx + sy * a.csy;
    int kf = 0;
    auto fetch = [&](uint32_t(&slot)[K]) {
        const buf_t r = buf_rsrc(fp, kf < len ? cellb : 0u);
        buf_load<K, 0, 2>(r, lane_off, slot); // read once: non-temporal
        fp += stepb;
        kf++;
    };
    uint32_t ring[PF][K];
#pragma unroll
    for (int u = 0; u < PF; u++) fetch(ring[u]);

    const buf_t pr = buf_rsrc(a.planes + (size_t)a.plane[s] * a.plane_bytes, (uint32_t)a.plane_bytes);
    uint32_t lo = (uint32_t)((((size_t)sy * a.W + sx) * a.D + d0) * 4);
    const uint32_t dl = (uint32_t)((dy * a.W + dx) * a.D * 4);
    const int nvalid = a.D - d0; // cells of this lane below D (may be <= 0)

    float Lp[K];
#pragma unroll
    for (int q = 0; q < K; q++) Lp[q] = O::inv();
    float m = O::inv(); // every cell of L' invalid: the next step starts the path

    auto body = [&](uint32_t(&slot)[K]) {
        float c[K];
#pragma unroll
        for (int q = 0; q < K; q++) c[q] = q < nvalid ? O::read(__builtin_bit_cast(float, slot[q])) : O::inv();
        fetch(slot);
        float Ln[K];
        if (!O::valid(m)) { // wave-uniform: path start or restart
#pragma unroll
            for (int q = 0; q < K; q++) Ln[q] = c[q];
        } else {
            const float left = wide_dpp<0x138>(Lp[K - 1]); // wave_shr:1
            const float right = wide_dpp<0x130>(Lp[0]);    // wave_shl:1
#pragma unroll
            for (int q = 0; q < K; q++) {
                const float lm = q == 0 ? left : Lp[q - 1];
                const float lp = q == K - 1 ? right : Lp[q + 1];
                const float t = O::mn(Lp[q], O::add(O::mn(lm, lp), a.P1));
                Ln[q] = O::add(c[q], O::excess(t, m, a.P2));
            }
        }
        float mn = Ln[0];
#pragma unroll
        for (int q = 1; q < K; q++) mn = O::mn(mn, Ln[q]);
        m = wave_min(mn);
        uint32_t w[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            Lp[q] = Ln[q];
            w[q] = __builtin_bit_cast(uint32_t, Ln[q]);
        }
        if (nvalid >= K) {
            buf_store<K>(pr, lo, w);
        } else {
#pragma unroll
            for (int q = 0; q < K; q++)
                if (q < nvalid) __builtin_amdgcn_raw_buffer_store_b32(w[q], pr, lo + 4 * q, 0, BUF_STORE_NT);
        }
        lo += dl;
    };
    int k0 = 0;
    for (; k0 + PF <= len; k0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) body(ring[u]);
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (k0 + u < len) body(ring[u]);
}

// S(x, y, d) = L_0 + L_1 + ... over the n planes in order (increasing
// direction index), one thread per pixel and 4 consecutive d; VEC: D % 4 ==
// 0 and 16-byte aligned sum cells.
template <bool VEC>
__global__ __launch_bounds__(256) void sgm_wide_fsum_kernel(const float *__restrict__ planes,
                                                            size_t plane_elems, int n,
                                                            uint8_t *__restrict__ S, int64_t ssx,
                                                            int64_t ssy, int W, int H, int D)
{
    const int G = (D + 3) / 4;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)W * H * G) return;
    const int g = (int)(idx % G);
    const int64_t pix = idx / G;
    const int x = (int)(pix % W), y = (int)(pix / W);
    const int d = 4 * g;
    const float *lp = planes + pix * D + d;
    float *sp = (float *)(S + x * ssx + y * ssy) + d;
    if constexpr (VEC) {
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        f32x4 acc = __builtin_nontemporal_load((const f32x4 *)lp);
        for (int r = 1; r < n; r++)
            acc = acc + __builtin_nontemporal_load((const f32x4 *)(lp + (size_t)r * plane_elems));
        *(f32x4 *)sp = acc;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (d + q >= D) break;
            float acc = lp[q];
            for (int r = 1; r < n; r++) acc = acc + lp[(size_t)r * plane_elems + q];
            sp[q] = acc;
        }
    }
}

// The float planes' sum + lowest-index WTA without writing S, as
// sgm_wide_sum_wta_kernel.
template <int P, bool VEC>
__global__ __launch_bounds__(256) void sgm_wide_fsum_wta_kernel(const float *__restrict__ planes,
                                                                size_t plane_elems, int n, int W, int H,
                                                                int D, int16_t *__restrict__ disp,
                                                                int pitch)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t pix = idx / P;
    const int t = (int)(idx % P);
    const bool live = pix < (int64_t)W * H; // every lane joins the shuffles
    const int x = live ? (int)(pix % W) : 0, y = live ? (int)(pix / W) : 0;
    const int d = 4 * t;
    uint32_t bv = 0xFFFFFFFFu, bd = 0xFFFFFFFFu;
    if (live && d < D) {
        const float *lp = planes + pix * D + d;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (VEC) {
            typedef float f32x4 __attribute__((ext_vector_type(4)));
            f32x4 v = __builtin_nontemporal_load((const f32x4 *)lp);
            for (int r = 1; r < n; r++)
                v = v + __builtin_nontemporal_load((const f32x4 *)(lp + (size_t)r * plane_elems));
            acc[0] = v.x, acc[1] = v.y, acc[2] = v.z, acc[3] = v.w;
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (d + q >= D) break;
                float a = lp[q];
                for (int r = 1; r < n; r++) a = a + lp[(size_t)r * plane_elems + q];
                acc[q] = a;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t k = wide_key(acc[q]);
            if (d + q < D && k < bv) {
                bv = k;
                bd = (uint32_t)(d + q);
            }
        }
    }
    const int best = pixel_argmin<P>(bv, bd);
    if (live && t == 0) disp[(size_t)y * pitch + x] = (int16_t)best;
}

template <int K>
static void launch_wide_l_k(const WideFArgs &a, hipStream_t s)
{
    constexpr int PF = 8;
    sgm_wide_l_kernel<K, PF><<<a.begin[a.ndirs], 64, 0, s>>>(a);
}

// The float-plane form for float volumes (planes below 2 GiB).
static int run_wide_f(tdvh_ctx *ctx, const tdvh_dsi *C, const tdvh_sgm_params *p, const tdvh_dsi *S,
                      int16_t *disp, int disp_pitch, hipStream_t s)
{
    const int W = C->width, H = C->height, D = C->depth;
    WideFArgs a{};
    a.C = (const uint8_t *)C->ptr;
    a.csx = C->stride_x * 4;
    a.csy = C->stride_y * 4;
    a.plane_bytes = (size_t)W * H * D * 4;
    a.W = W;
    a.H = H;
    a.D = D;
    a.P1 = p->fp1;
    a.P2 = p->fp2;
    // longest lines first: horizontal, vertical, then the diagonals
    int ord[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    if (W < H) {
        ord[0] = 2; ord[1] = 3; ord[2] = 0; ord[3] = 1;
    }
    a.begin[0] = 0;
    for (int i = 0; i < 8; i++) {
        const int dir = ord[i];
        if (!(p->dir_mask & (1 << dir))) continue;
        a.dir[a.ndirs] = dir;
        a.plane[a.ndirs] = popcount8(p->dir_mask & ((1 << dir) - 1));
        a.begin[a.ndirs + 1] = a.begin[a.ndirs] + dir_lines(dir, W, H);
        a.ndirs++;
    }
    if (a.ndirs == 0) return TDVH_OK;
    void *pb = nullptr;
    TDVH_TRY(ctx->scratch("sgm_wide_fplanes", a.plane_bytes * a.ndirs, &pb));
    a.planes = (uint8_t *)pb;
    prof_begin(s);
    switch ((D + 63) / 64) {
    case 1: launch_wide_l_k<1>(a, s); break;
    case 2: launch_wide_l_k<2>(a, s); break;
    case 3: launch_wide_l_k<3>(a, s); break;
    default: launch_wide_l_k<4>(a, s); break;
    }
    TDVH_TRY(check_launch("sgm_wide_lines_kernel", s));
    const size_t pe = (size_t)W * H * D;
    if (!S) { // disparities only
        const int P = sum_wta_lanes(D);
        const unsigned wblocks = (unsigned)(((int64_t)W * H * P + 255) / 256);
        prof_begin(s);
        switch (P) {
#define CASE(q)                                                                                    \
    case q:                                                                                        \
        if (D % 4 == 0) sgm_wide_fsum_wta_kernel<q, true><<<wblocks, 256, 0, s>>>((const float *)a.planes, pe, a.ndirs, W, H, D, disp, disp_pitch); \
        else sgm_wide_fsum_wta_kernel<q, false><<<wblocks, 256, 0, s>>>((const float *)a.planes, pe, a.ndirs, W, H, D, disp, disp_pitch); \
        break;
            CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
        }
        return check_launch("sgm_wide_sum_wta_kernel", s);
    }
    const int64_t threads = (int64_t)W * H * ((D + 3) / 4);
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    const bool vec = D % 4 == 0 && S->stride_x % 4 == 0 && S->stride_y % 4 == 0 &&
                     ((uintptr_t)S->ptr & 15) == 0;
    prof_begin(s);
    if (vec)
        sgm_wide_fsum_kernel<true><<<blocks, 256, 0, s>>>((const float *)a.planes, pe, a.ndirs,
                                                          (uint8_t *)S->ptr, S->stride_x * 4,
                                                          S->stride_y * 4, W, H, D);
    else
        sgm_wide_fsum_kernel<false><<<blocks, 256, 0, s>>>((const float *)a.planes, pe, a.ndirs,
                                                           (uint8_t *)S->ptr, S->stride_x * 4,
                                                           S->stride_y * 4, W, H, D);
    return check_launch("sgm_wide_sum_kernel", s);
}

template <typename T>
static int run_wide(const WideArgs<T> &a, int dir_mask, hipStream_t s)
{
    bool first = true;
    for (int dir = 0; dir < 8; dir++) {
        if (!(dir_mask & (1 << dir))) continue;
        prof_begin(s);
        switch ((a.D + 63) / 64) {
        case 1: launch_wide_k<T, 1>(a, dir, first, s); break;
        case 2: launch_wide_k<T, 2>(a, dir, first, s); break;
        case 3: launch_wide_k<T, 3>(a, dir, first, s); break;
        default: launch_wide_k<T, 4>(a, dir, first, s); break;
        }
        TDVH_TRY(check_launch("sgm_wide_dir_kernel", s));
        first = false;
    }
    return TDVH_OK;
}

// Float planes for the one-launch form: W*H*D floats addressed by 32-bit
// buffer offsets; TDVH_SGM_WIDE=dirs keeps the launch per direction.
static bool float_planes_fit(const tdvh_dsi *C)
{
    static const bool per_dir = [] {
        const char *e = getenv("TDVH_SGM_WIDE");
        return e && std::string(e) == "dirs";
    }();
    return !per_dir && (size_t)C->width * C->height * C->depth * 4 < ((size_t)1 << 31);
}

// The one-launch forms (byte planes for uint32 with P2 <= 255, float planes
// below 2 GiB), which can also feed the WTA from the sum without S.
bool sgm_wide_one_launch(const tdvh_dsi *C, const tdvh_sgm_params *p)
{
    if (C->dtype == TDVH_U32)
        return p->p2 <= 255 && (size_t)C->width * C->height * C->depth < ((size_t)1 << 31);
    return float_planes_fit(C);
}

int run_sgm_wide(tdvh_ctx *ctx, const tdvh_dsi *C, const tdvh_sgm_params *p, const tdvh_dsi *S,
                 int16_t *disp, int disp_pitch, hipStream_t s)
{
    TDVH_REQUIRE(S || sgm_wide_one_launch(C, p), "internal: S needed for this volume");
    if (C->dtype == TDVH_U32 && sgm_wide_one_launch(C, p))
        return run_wide_e(ctx, C, p, S, disp, disp_pitch, s);
    if (C->dtype == TDVH_U32) {
        WideArgs<uint32_t> a{(const uint32_t *)C->ptr, C->stride_x, C->stride_y,
                             (uint32_t *)S->ptr,       S->stride_x, S->stride_y,
                             C->width,                 C->height,   C->depth,
                             (uint32_t)p->p1,          (uint32_t)p->p2};
        return run_wide(a, p->dir_mask, s);
    }
    if (float_planes_fit(C)) return run_wide_f(ctx, C, p, S, disp, disp_pitch, s);
    WideArgs<float> a{(const float *)C->ptr, C->stride_x, C->stride_y, (float *)S->ptr,
                      S->stride_x,           S->stride_y, C->width,    C->height,
                      C->depth,              p->fp1,      p->fp2};
    return run_wide(a, p->dir_mask, s);
}

} // namespace tdvh
