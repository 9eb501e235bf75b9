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
// DPP shifts (the vacated end lane reads the invalid value); m is a DPP row
// minimum and four readlanes. Costs and the running sum stream through a
// prefetch ring of PF steps.
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

int run_sgm_wide(const tdvh_dsi *C, const tdvh_sgm_params *p, const tdvh_dsi *S, hipStream_t s)
{
    if (C->dtype == TDVH_U32) {
        WideArgs<uint32_t> a{(const uint32_t *)C->ptr, C->stride_x, C->stride_y,
                             (uint32_t *)S->ptr,       S->stride_x, S->stride_y,
                             C->width,                 C->height,   C->depth,
                             (uint32_t)p->p1,          (uint32_t)p->p2};
        return run_wide(a, p->dir_mask, s);
    }
    WideArgs<float> a{(const float *)C->ptr, C->stride_x, C->stride_y, (float *)S->ptr,
                      S->stride_x,           S->stride_y, C->width,    C->height,
                      C->depth,              p->fp1,      p->fp2};
    return run_wide(a, p->dir_mask, s);
}

} // namespace tdvh
