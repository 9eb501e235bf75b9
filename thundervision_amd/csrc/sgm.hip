// Semi-global matching, standard geometry, on gfx950.
//
// Recurrence (semiglobalcpu.cpp:87-125 with constant penalties):
//   L(p0, d) = C(p0, d)
//   L(p, d)  = C(p, d) + min(L'(d), L'(d-1) + P1, L'(d+1) + P1, m + P2) - m
// where L' is the previous pixel's vector on the path and m = min_d L'(d).
// Invalid costs (0xFFFF) give invalid L; invalid operands drop out of the
// minimum (saturating uint16 adds keep 0xFFFF + P1 at 0xFFFF).
//
// Mapping: one 16-lane DPP row owns one scan line and holds its disparity
// vector in m packed uint16x2 registers per lane (SPLIT16 order: word j*m+r
// of lane j is d = j*m+r in the low half and d = 16m + j*m+r in the high
// half). A wave scans 4 lines; the d-1 / d+1 neighbours are register
// renames except at the lane boundary (one row_ror DPP move each way plus a
// fix-up on the end lanes), and m is a 4-step DPP row all-reduce. Every
// direction's lines are independent, so all 8 directions run in one launch
// and each writes its own path-cost plane; a second kernel sums the planes
// and takes the lowest-index argmin (wtacpu.cpp:20-33).
#include <cstdlib>

#include "kernels.hpp"

namespace tdvh {

struct LineSet {
    int ndirs;
    int dir[8];
    int wave_begin[9]; // prefix sums of waves (4 lines per wave) per slot
};

// Start pixel and length of standard-geometry line `line` of direction `dir`.
__device__ __forceinline__ void line_geom(int dir, int line, int W, int H, int &sx, int &sy,
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
    const uint32_t *C;  // SPLIT16 cost volume (staged), or null
    const uint32_t *LT; // 3 planes W*H: splatted left BT terms l2, lmin, lmax (fused)
    const uint32_t *RT; // 3 planes H*Wr: paired right BT terms r2, rmin, rmax (fused)
    int Wr;             // row length of RT: W + 16m
    uint8_t *planes;    // path-cost planes (E bytes or L words)
    size_t plane_bytes;
};

// Packed uint16 saturating subtract / max.
__device__ __forceinline__ uint32_t pk_subs(uint32_t a, uint32_t b)
{
    return u32(__builtin_elementwise_sub_sat(v2(a), v2(b)));
}
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b)
{
    return u32(__builtin_elementwise_max(v2(a), v2(b)));
}

// Invalid-cell mask of lane j, word r: d >= D (padding) or d > x.
template <int M>
__device__ __forceinline__ uint32_t cell_mask(int j, int r, int D, int x)
{
    const int dlo = j * M + r, dhi = 16 * M + j * M + r;
    return ((dlo >= D || dlo > x) ? 0x0000FFFFu : 0u) | ((dhi >= D || dhi > x) ? 0xFFFF0000u : 0u);
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

// One SGM step on the packed vector of a line. Lp: previous L (in/out),
// c: costs of the current pixel, m: previous minimum (in/out); E receives
// L - C = min(...) - m, in [0, P2] for valid cells.
template <int M>
__device__ __forceinline__ void sgm_step(uint32_t (&Lp)[M], const uint32_t (&c)[M], uint32_t &m,
                                         int j, uint32_t p1, uint32_t P2, uint32_t (&E)[M])
{
    uint32_t A = row_ror1(Lp[M - 1]);
    if (j == 0) A = (A << 16) | 0xFFFFu; // d-1 of d=0 is absent; of d=16m is lane 15's low
    uint32_t B = row_ror15(Lp[0]);
    if (j == 15) B = (B >> 16) | 0xFFFF0000u; // d+1 of d=16m-1 is lane 0's high; of Dp-1 absent
    const uint32_t mm = pk_splat(m), mp2 = pk_splat(m + P2);
    uint32_t nmin = 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < M; r++) {
        const uint32_t dm1 = r == 0 ? A : Lp[r - 1];
        const uint32_t dp1 = r == M - 1 ? B : Lp[r + 1];
        uint32_t t = pk_adds(pk_min(dm1, dp1), p1);
        t = pk_min(pk_min(t, Lp[r]), mp2);
        E[r] = pk_sub(t, mm);
    }
#pragma unroll
    for (int r = 0; r < M; r++) {
        Lp[r] = pk_adds(E[r], c[r]);
        nmin = pk_min(nmin, Lp[r]);
    }
    m = row_min_u32(min(nmin & 0xFFFFu, nmin >> 16));
}

// Low bytes of the two halves of words a (-> bytes 0,1) and b (-> bytes 2,3).
__device__ __forceinline__ uint32_t pack_e(uint32_t a, uint32_t b)
{
    return __builtin_amdgcn_perm(b, a, 0x06040200u);
}

template <int M, bool FUSED, bool EOUT, int PF>
__global__ __launch_bounds__(256) void sgm_lines_kernel(const SgmArgs a, const LineSet ls)
{
    constexpr int WPP = 16 * M;
    constexpr int NW = FUSED ? 3 + 3 * M : M; // prefetched words per step
    const int lane = threadIdx.x & 63;
    const int gw = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gw >= ls.wave_begin[ls.ndirs]) return;
    int s = 0;
    while (s < ls.ndirs - 1 && gw >= ls.wave_begin[s + 1]) s++;
    const int dir = ls.dir[s];
    const int line = (gw - ls.wave_begin[s]) * 4 + (lane >> 4);
    const int j = lane & 15;
    const int W = a.W, H = a.H;
    if (line >= dir_lines(dir, W, H)) return; // whole 16-lane row leaves together
    int sx, sy, len;
    line_geom(dir, line, W, H, sx, sy, len);
    const int dx = dir_dx(dir), dy = dir_dy(dir);
    const long step = (long)dy * W + dx;
    const long p0 = (long)sy * W + sx;
    uint8_t *outp = a.planes + (size_t)s * a.plane_bytes;
    const size_t lt_plane = (size_t)W * H, rt_plane = (size_t)a.Wr * H;

    uint32_t ring[PF][NW];
    auto fetch = [&](uint32_t(&slot)[NW], int k) {
        const int kk = k < len ? k : len - 1;
        const long p = p0 + kk * step;
        if constexpr (!FUSED) {
            const uint32_t *src = a.C + (size_t)p * WPP + j * M;
#pragma unroll
            for (int r = 0; r < M; r++) slot[r] = src[r];
        } else {
            const int x = sx + kk * dx, y = sy + kk * dy;
#pragma unroll
            for (int t = 0; t < 3; t++) slot[t] = a.LT[t * lt_plane + p];
            const uint32_t *rt = a.RT + (size_t)y * a.Wr + (W - 1 - x) + j * M;
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int r = 0; r < M; r++) slot[3 + t * M + r] = rt[t * rt_plane + r];
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

    for (int k0 = 0; k0 < len; k0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int k = k0 + u;
            if (k < len) {
                uint32_t c[M];
                if constexpr (!FUSED) {
#pragma unroll
                    for (int r = 0; r < M; r++) c[r] = ring[u][r] | padm[r];
                } else {
                    const uint32_t lt[3] = {ring[u][0], ring[u][1], ring[u][2]};
                    uint32_t r2[M], rmn[M], rmx[M];
#pragma unroll
                    for (int r = 0; r < M; r++) {
                        r2[r] = ring[u][3 + r];
                        rmn[r] = ring[u][3 + M + r];
                        rmx[r] = ring[u][3 + 2 * M + r];
                    }
                    bt_costs<M>(lt, r2, rmn, rmx, j, a.D, sx + k * dx, c);
                }
                fetch(ring[u], k + PF);
                uint32_t E[M];
                sgm_step<M>(Lp, c, m, j, p1, (uint32_t)a.P2, E);
                const size_t p = (size_t)(p0 + k * step);
                if constexpr (EOUT) {
                    // E in bytes: lane j holds 2M bytes at (p*32M + j*2M)
                    uint8_t *dst = outp + p * (32 * M) + j * (2 * M);
                    if constexpr (M % 2 == 0) {
#pragma unroll
                        for (int r = 0; r < M; r += 2)
                            ((uint32_t *)dst)[r / 2] = pack_e(E[r], E[r + 1]);
                    } else {
#pragma unroll
                        for (int r = 0; r < M; r++)
                            ((uint16_t *)dst)[r] = (uint16_t)pack_e(E[r], 0);
                    }
                } else {
                    uint32_t *dst = (uint32_t *)outp + p * WPP + j * M;
#pragma unroll
                    for (int r = 0; r < M; r++) dst[r] = Lp[r];
                }
            }
        }
    }
}

int popcount8(int mask)
{
    int n = 0;
    for (int i = 0; i < 8; i++) n += (mask >> i) & 1;
    return n;
}

// Prefetch depth (steps of cost data in flight per line): 4, 8 or 16.
static int lines_prefetch()
{
    static int pf = [] {
        const char *e = getenv("TDVH_SGM_PREFETCH");
        const int v = e ? atoi(e) : 4;
        return (v == 8 || v == 16) ? v : 4;
    }();
    return pf;
}

template <int M, bool FUSED, bool EOUT>
static void launch_lines_k(const SgmArgs &a, const LineSet &ls, hipStream_t s)
{
    const int waves = ls.wave_begin[ls.ndirs];
    const int blocks = (waves + 3) / 4;
    switch (lines_prefetch()) {
    case 8: sgm_lines_kernel<M, FUSED, EOUT, 8><<<blocks, 256, 0, s>>>(a, ls); break;
    case 16: sgm_lines_kernel<M, FUSED, EOUT, 16><<<blocks, 256, 0, s>>>(a, ls); break;
    default: sgm_lines_kernel<M, FUSED, EOUT, 4><<<blocks, 256, 0, s>>>(a, ls); break;
    }
}

template <int M>
static void launch_lines_m(const SgmArgs &a, const LineSet &ls, bool eout, hipStream_t s)
{
    const bool fused = a.C == nullptr;
    if (fused && eout) launch_lines_k<M, true, true>(a, ls, s);
    else if (fused) launch_lines_k<M, true, false>(a, ls, s);
    else if (eout) launch_lines_k<M, false, true>(a, ls, s);
    else launch_lines_k<M, false, false>(a, ls, s);
}

static LineSet make_lineset(int W, int H, int dir_mask)
{
    LineSet ls{};
    ls.wave_begin[0] = 0;
    for (int dir = 0; dir < 8; dir++) {
        if (!(dir_mask & (1 << dir))) continue;
        ls.dir[ls.ndirs] = dir;
        ls.wave_begin[ls.ndirs + 1] = ls.wave_begin[ls.ndirs] + (dir_lines(dir, W, H) + 3) / 4;
        ls.ndirs++;
    }
    return ls;
}

int launch_sgm_lines(const SgmArgs &a, int dir_mask, bool eout, hipStream_t s)
{
    const LineSet ls = make_lineset(a.W, a.H, dir_mask);
    if (ls.ndirs == 0) return TDVH_OK;
    prof_begin(s);
    switch (split_m(a.D)) {
#define CASE(m)                                                                                    \
    case m: launch_lines_m<m>(a, ls, eout, s); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: set_error("depth %d above %d", a.D, MAX_DEPTH); return TDVH_ERR_UNSUPPORTED;
    }
    return check_launch("sgm_lines_kernel", s);
}

// ---- BT term planes for the fused path ---------------------------------------

// LT[t][y][x]: splatted left terms; RT[t][y][i]: right terms of pixels
// W-1-i (low half) and W-1-i-16m (high half), clamped at x' = 0.
__global__ __launch_bounds__(256) void bt_planes_kernel(const uint8_t *__restrict__ L,
                                                        const uint8_t *__restrict__ R, int W,
                                                        int H, int pitch, int m,
                                                        uint32_t *__restrict__ LT,
                                                        uint32_t *__restrict__ RT, int Wr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const uint8_t *lr = L + (size_t)y * pitch;
    const uint8_t *rr = R + (size_t)y * pitch;
    auto terms = [&](const uint8_t *row, int x, int &v2, int &mn, int &mx) {
        x = x < 0 ? 0 : x;
        const int c = row[x];
        const int a = c + row[x > 0 ? x - 1 : 0], b = c + row[x < W - 1 ? x + 1 : W - 1];
        v2 = 2 * c;
        mn = min(v2, min(a, b));
        mx = max(v2, max(a, b));
    };
    if (i < W) {
        int v2, mn, mx;
        terms(lr, i, v2, mn, mx);
        const size_t o = (size_t)y * W + i, pl = (size_t)W * H;
        LT[o] = pk_splat(v2);
        LT[pl + o] = pk_splat(mn);
        LT[2 * pl + o] = pk_splat(mx);
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

int launch_bt_planes(const uint8_t *L, const uint8_t *R, int W, int H, int pitch, int D,
                     uint32_t *LT, uint32_t *RT, hipStream_t s)
{
    const int m = split_m(D), Wr = W + 16 * m;
    dim3 grid((Wr + 255) / 256, H);
    prof_begin(s);
    bt_planes_kernel<<<grid, 256, 0, s>>>(L, R, W, H, pitch, m, LT, RT, Wr);
    return check_launch("bt_planes_kernel", s);
}

// ---- plane sum + lowest-index WTA ----------------------------------------

// Unpacks the two bytes of half h of word w into a packed uint16x2.
__device__ __forceinline__ uint32_t unpack_e(uint32_t w, int h)
{
    return __builtin_amdgcn_perm(0u, w, h ? 0x0C030C02u : 0x0C010C00u);
}

template <int M, bool FUSED, bool EOUT>
__global__ __launch_bounds__(256) void sgm_sum_wta_kernel(const SgmArgs a, int nplanes,
                                                          uint32_t *__restrict__ S,
                                                          int16_t *__restrict__ disp, int pitch)
{
    constexpr int WPP = 16 * M;
    const int W = a.W, H = a.H;
    const long p = (long)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int j = threadIdx.x & 15;
    if (p >= (long)W * H) return;
    const int y = (int)(p / W), x = (int)(p - (long)y * W);
    uint32_t acc[M];
    if constexpr (EOUT) {
        // S = n*C + sum of E (invalid where C is)
        uint32_t c[M];
        if constexpr (FUSED) {
            const size_t lt_plane = (size_t)W * H, rt_plane = (size_t)a.Wr * H;
            const uint32_t lt[3] = {a.LT[p], a.LT[lt_plane + p], a.LT[2 * lt_plane + p]};
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

template <int M>
static void launch_sum_m(const SgmArgs &a, int nplanes, bool eout, uint32_t *S, int16_t *disp,
                         int pitch, int blocks, hipStream_t s)
{
    const bool fused = a.C == nullptr;
    if (eout && fused)
        sgm_sum_wta_kernel<M, true, true><<<blocks, 256, 0, s>>>(a, nplanes, S, disp, pitch);
    else if (eout)
        sgm_sum_wta_kernel<M, false, true><<<blocks, 256, 0, s>>>(a, nplanes, S, disp, pitch);
    else
        sgm_sum_wta_kernel<M, false, false><<<blocks, 256, 0, s>>>(a, nplanes, S, disp, pitch);
}

int launch_sgm_sum_wta(const SgmArgs &a, int nplanes, bool eout, uint32_t *S_split, int16_t *disp,
                       int disp_pitch, hipStream_t s)
{
    const long pixels = (long)a.W * a.H;
    const int blocks = (int)((pixels + 15) / 16);
    prof_begin(s);
    switch (split_m(a.D)) {
#define CASE(m)                                                                                    \
    case m: launch_sum_m<m>(a, nplanes, eout, S_split, disp, disp_pitch, blocks, s); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: set_error("depth %d above %d", a.D, MAX_DEPTH); return TDVH_ERR_UNSUPPORTED;
    }
    return check_launch("sgm_sum_wta_kernel", s);
}

} // namespace tdvh

using namespace tdvh;

static int validate_sgm_params(const tdvh_sgm_params *p)
{
    TDVH_REQUIRE(p != nullptr, "params is null");
    TDVH_REQUIRE(p->p1 >= 0 && p->p2 >= 0 && p->p1 <= 1024 && p->p2 <= 1024,
                 "penalties out of range (P1=%d, P2=%d)", p->p1, p->p2);
    TDVH_REQUIRE((p->dir_mask & 0xFF) != 0 && (p->dir_mask & ~0xFF) == 0, "bad dir_mask 0x%x",
                 p->dir_mask);
    return TDVH_OK;
}

// Standard-geometry pipeline from a SPLIT16 cost volume (Csplit) or from
// the images (Csplit null: BT recomputed from term planes, no C volume).
static int sgm_std(tdvh_ctx *ctx, const uint32_t *Csplit, const uint8_t *left,
                   const uint8_t *right, int pitch, int W, int H, int D,
                   const tdvh_sgm_params *params, tdvh_dsi *S_out, int16_t *disp, int disp_pitch,
                   hipStream_t s)
{
    const int nplanes = popcount8(params->dir_mask);
    const int m = split_m(D);
    // path-cost planes: L - C in bytes when P2 fits a byte, else L words
    const bool eout = params->p2 <= 255;
    SgmArgs a{};
    a.W = W;
    a.H = H;
    a.D = D;
    a.P1 = params->p1;
    a.P2 = params->p2;
    a.C = Csplit;
    a.Wr = W + 16 * m;
    a.plane_bytes = (size_t)W * H * 32 * m * (eout ? 1 : 2);
    void *pb = nullptr;
    TDVH_TRY(ctx->scratch("sgm_planes", a.plane_bytes * nplanes, &pb));
    a.planes = (uint8_t *)pb;
    if (!Csplit) {
        void *tb = nullptr;
        const size_t words = 3 * (size_t)W * H + 3 * (size_t)a.Wr * H;
        TDVH_TRY(ctx->scratch("bt_planes", words * 4, &tb));
        a.LT = (const uint32_t *)tb;
        a.RT = (const uint32_t *)tb + 3 * (size_t)W * H;
        TDVH_TRY(launch_bt_planes(left, right, W, H, pitch, D, (uint32_t *)a.LT,
                                  (uint32_t *)a.RT, s));
        if (!eout) {
            // L words need C in the sum: fall back to a staged volume
            void *cb = nullptr;
            TDVH_TRY(ctx->scratch("cost_c", tdvh_dsi_bytes(W, H, D, TDVH_U16, TDVH_LAYOUT_SPLIT16),
                                  &cb));
            void *tl = nullptr;
            TDVH_TRY(ctx->scratch("bt_terms", (size_t)W * H * 8, &tl));
            uint32_t *TL = (uint32_t *)tl, *TR = TL + (size_t)W * H;
            TDVH_TRY(launch_bt_terms(left, right, W, H, pitch, TL, TR, s));
            TDVH_TRY(launch_bt_split(TL, TR, W, H, D, (uint32_t *)cb, s));
            a.C = (const uint32_t *)cb;
        }
    }
    TDVH_TRY(launch_sgm_lines(a, params->dir_mask, eout, s));
    uint32_t *Ssplit = nullptr;
    if (S_out) {
        TDVH_REQUIRE(S_out->width == W && S_out->height == H && S_out->depth == D &&
                         S_out->dtype == TDVH_U16 && S_out->ptr,
                     "S_out must be a uint16 %dx%dx%d volume", W, H, D);
        if (S_out->layout == TDVH_LAYOUT_SPLIT16)
            Ssplit = (uint32_t *)S_out->ptr;
        else {
            void *sb = nullptr;
            TDVH_TRY(ctx->scratch("sgm_s", (size_t)W * H * 64 * m, &sb));
            Ssplit = (uint32_t *)sb;
        }
    }
    TDVH_TRY(launch_sgm_sum_wta(a, nplanes, eout, Ssplit, disp, disp_pitch, s));
    if (S_out && S_out->layout != TDVH_LAYOUT_SPLIT16)
        TDVH_TRY(launch_convert((uint16_t *)S_out->ptr, S_out->stride_x, S_out->stride_y, Ssplit,
                                W, H, D, false, s));
    return TDVH_OK;
}

extern "C" int tdvh_sgm(tdvh_ctx *ctx, const tdvh_dsi *C, const tdvh_sgm_params *params,
                        tdvh_dsi *S_out, int16_t *disp, int disp_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_TRY(validate_sgm_params(params));
    TDVH_REQUIRE(C && C->ptr && disp, "null pointer argument");
    TDVH_REQUIRE(C->dtype == TDVH_U16, "SGM needs a uint16 cost volume");
    const int W = C->width, H = C->height, D = C->depth;
    TDVH_REQUIRE(W > 0 && H > 0 && D > 0 && D <= MAX_DEPTH, "bad volume %dx%dx%d (D <= %d)", W, H,
                 D, MAX_DEPTH);
    TDVH_REQUIRE(disp_pitch >= W, "disp_pitch %d < width %d", disp_pitch, W);
    hipStream_t s = (hipStream_t)stream;
    const uint32_t *Csplit = (const uint32_t *)C->ptr;
    if (C->layout != TDVH_LAYOUT_SPLIT16) {
        void *cb = nullptr;
        TDVH_TRY(ctx->scratch("sgm_c", tdvh_dsi_bytes(W, H, D, TDVH_U16, TDVH_LAYOUT_SPLIT16), &cb));
        TDVH_TRY(launch_convert((uint16_t *)C->ptr, C->stride_x, C->stride_y, (uint32_t *)cb, W, H,
                                D, true, s));
        Csplit = (const uint32_t *)cb;
    }
    if (params->flags & TDVH_SGM_REF_GEOMETRY) {
        TDVH_REQUIRE(D >= 2, "reference geometry needs D >= 2");
        int32_t *Sref = nullptr;
        if (S_out) {
            TDVH_REQUIRE(S_out->width == W && S_out->height == H && S_out->depth == D &&
                             S_out->dtype == TDVH_I32 && S_out->layout == TDVH_LAYOUT_STRIDED &&
                             S_out->stride_x == D && S_out->stride_y == (int64_t)W * D &&
                             S_out->ptr,
                         "S_out must be a row-major int32 %dx%dx%d volume", W, H, D);
            Sref = (int32_t *)S_out->ptr;
        } else {
            void *sb = nullptr;
            TDVH_TRY(ctx->scratch("sgm_sref", (size_t)W * H * D * 4, &sb));
            Sref = (int32_t *)sb;
        }
        TDVH_TRY(run_sgm_ref(ctx, Csplit, W, H, D, params->p1, params->p2, params->flags, Sref, s));
        tdvh_dsi sd;
        TDVH_TRY(tdvh_dsi_init(&sd, Sref, W, H, D, TDVH_I32, TDVH_LAYOUT_STRIDED));
        return tdvh_wta(ctx, &sd, TDVH_WTA_LOWEST, disp, disp_pitch, stream);
    }
    return sgm_std(ctx, Csplit, nullptr, nullptr, 0, W, H, D, params, S_out, disp, disp_pitch, s);
}

extern "C" int tdvh_stereo_bt_sgm(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right,
                                  int width, int height, int pitch, int depth,
                                  const tdvh_sgm_params *params, int fusion, int16_t *disp,
                                  int disp_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_TRY(validate_sgm_params(params));
    TDVH_REQUIRE(left && right && disp, "null pointer argument");
    TDVH_REQUIRE(width > 0 && height > 0 && pitch >= width, "bad image %dx%d pitch %d", width,
                 height, pitch);
    TDVH_REQUIRE(depth > 0 && depth <= MAX_DEPTH, "depth %d outside [1, %d]", depth, MAX_DEPTH);
    TDVH_REQUIRE(disp_pitch >= width, "disp_pitch %d < width %d", disp_pitch, width);
    TDVH_REQUIRE(!(params->flags & TDVH_SGM_REF_GEOMETRY) || fusion == TDVH_FUSE_STAGED,
                 "reference geometry runs staged only");
    hipStream_t s = (hipStream_t)stream;
    if (fusion == TDVH_FUSE_COST)
        return sgm_std(ctx, nullptr, left, right, pitch, width, height, depth, params, nullptr,
                       disp, disp_pitch, s);
    TDVH_REQUIRE(fusion == TDVH_FUSE_STAGED, "unknown fusion mode %d", fusion);
    void *tl = nullptr;
    TDVH_TRY(ctx->scratch("bt_terms", (size_t)width * height * 8, &tl));
    uint32_t *TL = (uint32_t *)tl, *TR = TL + (size_t)width * height;
    TDVH_TRY(launch_bt_terms(left, right, width, height, pitch, TL, TR, s));
    void *cb = nullptr;
    TDVH_TRY(ctx->scratch("cost_c", tdvh_dsi_bytes(width, height, depth, TDVH_U16,
                                                   TDVH_LAYOUT_SPLIT16),
                          &cb));
    TDVH_TRY(launch_bt_split(TL, TR, width, height, depth, (uint32_t *)cb, s));
    tdvh_dsi cd;
    TDVH_TRY(tdvh_dsi_init(&cd, cb, width, height, depth, TDVH_U16, TDVH_LAYOUT_SPLIT16));
    return tdvh_sgm(ctx, &cd, params, nullptr, disp, disp_pitch, stream);
}
