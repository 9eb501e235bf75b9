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

// One SGM step on the packed vector of a line. Lp: previous L (in/out),
// c: costs of the current pixel, m: previous minimum (in/out).
template <int M>
__device__ __forceinline__ void sgm_step(uint32_t (&Lp)[M], const uint32_t (&c)[M], uint32_t &m,
                                         int j, uint32_t p1, uint32_t P2)
{
    uint32_t A = row_ror1(Lp[M - 1]);
    if (j == 0) A = (A << 16) | 0xFFFFu; // d-1 of d=0 is absent; of d=16m is lane 15's low
    uint32_t B = row_ror15(Lp[0]);
    if (j == 15) B = (B >> 16) | 0xFFFF0000u; // d+1 of d=16m-1 is lane 0's high; of Dp-1 absent
    const uint32_t mm = pk_splat(m), mp2 = pk_splat(m + P2);
    uint32_t Ln[M];
    uint32_t nmin = 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < M; r++) {
        const uint32_t dm1 = r == 0 ? A : Lp[r - 1];
        const uint32_t dp1 = r == M - 1 ? B : Lp[r + 1];
        uint32_t t = pk_adds(pk_min(dm1, dp1), p1);
        t = pk_min(pk_min(t, Lp[r]), mp2);
        Ln[r] = pk_adds(pk_sub(t, mm), c[r]);
        nmin = pk_min(nmin, Ln[r]);
    }
#pragma unroll
    for (int r = 0; r < M; r++) Lp[r] = Ln[r];
    m = row_min_u32(min(nmin & 0xFFFFu, nmin >> 16));
}

template <int M, bool FUSED, int PF>
__global__ __launch_bounds__(256) void sgm_lines_kernel(
    const uint32_t *__restrict__ C, const uint32_t *__restrict__ TL,
    const uint32_t *__restrict__ TR, int W, int H, int D, int P1, int P2,
    uint32_t *__restrict__ Lout, size_t plane_words, LineSet ls)
{
    constexpr int WPP = 16 * M;
    constexpr int NW = FUSED ? 2 * M + 1 : M; // prefetched words per step
    const int lane = threadIdx.x & 63;
    const int gw = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gw >= ls.wave_begin[ls.ndirs]) return;
    int s = 0;
    while (s < ls.ndirs - 1 && gw >= ls.wave_begin[s + 1]) s++;
    const int dir = ls.dir[s];
    const int line = (gw - ls.wave_begin[s]) * 4 + (lane >> 4);
    const int j = lane & 15;
    if (line >= dir_lines(dir, W, H)) return; // whole 16-lane row leaves together
    int sx, sy, len;
    line_geom(dir, line, W, H, sx, sy, len);
    const int dx = dir_dx(dir);
    const long step = (long)dir_dy(dir) * W + dx;
    const long p0 = (long)sy * W + sx;
    uint32_t *out = Lout + (size_t)s * plane_words + j * M;

    uint32_t ring[PF][NW];
    auto fetch = [&](uint32_t(&slot)[NW], int k) {
        const int kk = k < len ? k : len - 1;
        const long p = p0 + kk * step;
        if constexpr (!FUSED) {
            const uint32_t *src = C + (size_t)p * WPP + j * M;
#pragma unroll
            for (int r = 0; r < M; r++) slot[r] = src[r];
        } else {
            const int x = sx + kk * dx;
            slot[0] = TL[p];
#pragma unroll
            for (int r = 0; r < M; r++) {
                const int dlo = j * M + r, dhi = 16 * M + j * M + r;
                slot[1 + 2 * r] = TR[p - min(dlo, x)];
                slot[2 + 2 * r] = TR[p - min(dhi, x)];
            }
        }
    };
    auto costs = [&](const uint32_t(&slot)[NW], int k, uint32_t(&c)[M]) {
        if constexpr (!FUSED) {
#pragma unroll
            for (int r = 0; r < M; r++) c[r] = slot[r];
        } else {
            const int x = sx + k * dx;
#pragma unroll
            for (int r = 0; r < M; r++) {
                const int dlo = j * M + r, dhi = 16 * M + j * M + r;
                const uint32_t lo = (dlo < D && dlo <= x) ? bt_cost(slot[0], slot[1 + 2 * r]) : INV16;
                const uint32_t hi = (dhi < D && dhi <= x) ? bt_cost(slot[0], slot[2 + 2 * r]) : INV16;
                c[r] = lo | (hi << 16);
            }
        }
    };

#pragma unroll
    for (int u = 0; u < PF; u++) fetch(ring[u], u);

    // L' = 0 and m = 0 make the first step return L = C (path start).
    uint32_t Lp[M];
#pragma unroll
    for (int r = 0; r < M; r++) Lp[r] = 0;
    uint32_t m = 0;
    const uint32_t p1 = pk_splat(P1);

    for (int k0 = 0; k0 < len; k0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int k = k0 + u;
            if (k < len) {
                uint32_t c[M];
                costs(ring[u], k, c);
                fetch(ring[u], k + PF);
                sgm_step<M>(Lp, c, m, j, p1, (uint32_t)P2);
                uint32_t *dst = out + (size_t)(p0 + k * step) * WPP;
#pragma unroll
                for (int r = 0; r < M; r++) dst[r] = Lp[r];
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

template <int M>
static void launch_lines_m(const uint32_t *C, const uint32_t *TL, const uint32_t *TR, int W, int H,
                           int D, int P1, int P2, uint32_t *Lout, size_t plane, const LineSet &ls,
                           hipStream_t s)
{
    const int waves = ls.wave_begin[ls.ndirs];
    const int blocks = (waves + 3) / 4;
    if (TL)
        sgm_lines_kernel<M, true, 4><<<blocks, 256, 0, s>>>(C, TL, TR, W, H, D, P1, P2, Lout,
                                                            plane, ls);
    else
        sgm_lines_kernel<M, false, 4><<<blocks, 256, 0, s>>>(C, TL, TR, W, H, D, P1, P2, Lout,
                                                             plane, ls);
}

int launch_sgm_lines(const uint32_t *C, const uint32_t *TL, const uint32_t *TR, int W, int H,
                     int D, int P1, int P2, int dir_mask, uint32_t *Lout, size_t plane_words,
                     hipStream_t s)
{
    LineSet ls{};
    ls.wave_begin[0] = 0;
    for (int dir = 0; dir < 8; dir++) {
        if (!(dir_mask & (1 << dir))) continue;
        ls.dir[ls.ndirs] = dir;
        ls.wave_begin[ls.ndirs + 1] = ls.wave_begin[ls.ndirs] + (dir_lines(dir, W, H) + 3) / 4;
        ls.ndirs++;
    }
    if (ls.ndirs == 0) return TDVH_OK;
    prof_begin(s);
    switch (split_m(D)) {
#define CASE(m)                                                                                    \
    case m: launch_lines_m<m>(C, TL, TR, W, H, D, P1, P2, Lout, plane_words, ls, s); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: set_error("depth %d above %d", D, MAX_DEPTH); return TDVH_ERR_UNSUPPORTED;
    }
    return check_launch("sgm_lines_kernel", s);
}

// ---- plane sum + lowest-index WTA ----------------------------------------

template <int M>
__global__ __launch_bounds__(256) void sgm_sum_wta_kernel(const uint32_t *__restrict__ L,
                                                          size_t plane, int nplanes, int W, int H,
                                                          uint32_t *__restrict__ S,
                                                          int16_t *__restrict__ disp, int pitch)
{
    constexpr int WPP = 16 * M;
    const long p = (long)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int j = threadIdx.x & 15;
    if (p >= (long)W * H) return;
    const size_t base = (size_t)p * WPP + j * M;
    uint32_t acc[M];
#pragma unroll
    for (int r = 0; r < M; r++) acc[r] = L[base + r];
    for (int q = 1; q < nplanes; q++) {
        const uint32_t *src = L + q * plane + base;
#pragma unroll
        for (int r = 0; r < M; r++) acc[r] = pk_adds(acc[r], src[r]);
    }
    if (S) {
#pragma unroll
        for (int r = 0; r < M; r++) S[base + r] = acc[r];
    }
    uint32_t key = 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < M; r++) {
        const uint32_t dlo = j * M + r, dhi = 16 * M + j * M + r;
        key = min(key, ((acc[r] & 0xFFFFu) << 16) | dlo);
        key = min(key, (acc[r] & 0xFFFF0000u) | dhi);
    }
    key = row_min_u32(key);
    if (j == 0) {
        const int y = (int)(p / W), x = (int)(p - (long)y * W);
        disp[(size_t)y * pitch + x] = (int16_t)(key & 0xFFFFu);
    }
}

int launch_sgm_sum_wta(const uint32_t *L, size_t plane_words, int nplanes, int W, int H, int D,
                       uint32_t *S_split, int16_t *disp, int disp_pitch, hipStream_t s)
{
    const long pixels = (long)W * H;
    const int blocks = (int)((pixels + 15) / 16);
    prof_begin(s);
    switch (split_m(D)) {
#define CASE(m)                                                                                    \
    case m:                                                                                        \
        sgm_sum_wta_kernel<m><<<blocks, 256, 0, s>>>(L, plane_words, nplanes, W, H, S_split, disp, \
                                                     disp_pitch);                                  \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: set_error("depth %d above %d", D, MAX_DEPTH); return TDVH_ERR_UNSUPPORTED;
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

// Standard-geometry pipeline from a SPLIT16 cost volume or BT terms.
static int sgm_std(tdvh_ctx *ctx, const uint32_t *Csplit, const uint32_t *TL, const uint32_t *TR,
                   int W, int H, int D, const tdvh_sgm_params *params, tdvh_dsi *S_out,
                   int16_t *disp, int disp_pitch, hipStream_t s)
{
    const int nplanes = popcount8(params->dir_mask);
    const size_t plane = (size_t)W * H * split_words(D);
    void *lbuf = nullptr;
    TDVH_TRY(ctx->scratch("sgm_planes", plane * 4 * nplanes, &lbuf));
    TDVH_TRY(launch_sgm_lines(Csplit, TL, TR, W, H, D, params->p1, params->p2, params->dir_mask,
                              (uint32_t *)lbuf, plane, s));
    uint32_t *Ssplit = nullptr;
    if (S_out) {
        TDVH_REQUIRE(S_out->width == W && S_out->height == H && S_out->depth == D &&
                         S_out->dtype == TDVH_U16 && S_out->ptr,
                     "S_out must be a uint16 %dx%dx%d volume", W, H, D);
        if (S_out->layout == TDVH_LAYOUT_SPLIT16)
            Ssplit = (uint32_t *)S_out->ptr;
        else {
            void *sb = nullptr;
            TDVH_TRY(ctx->scratch("sgm_s", plane * 4, &sb));
            Ssplit = (uint32_t *)sb;
        }
    }
    TDVH_TRY(launch_sgm_sum_wta((const uint32_t *)lbuf, plane, nplanes, W, H, D, Ssplit, disp,
                                disp_pitch, s));
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
    return sgm_std(ctx, Csplit, nullptr, nullptr, W, H, D, params, S_out, disp, disp_pitch, s);
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
    void *tl = nullptr;
    TDVH_TRY(ctx->scratch("bt_terms", (size_t)width * height * 8, &tl));
    uint32_t *TL = (uint32_t *)tl, *TR = TL + (size_t)width * height;
    TDVH_TRY(launch_bt_terms(left, right, width, height, pitch, TL, TR, s));
    if (fusion == TDVH_FUSE_COST)
        return sgm_std(ctx, nullptr, TL, TR, width, height, depth, params, nullptr, disp,
                       disp_pitch, s);
    TDVH_REQUIRE(fusion == TDVH_FUSE_STAGED, "unknown fusion mode %d", fusion);
    void *cb = nullptr;
    TDVH_TRY(ctx->scratch("cost_c", tdvh_dsi_bytes(width, height, depth, TDVH_U16,
                                                   TDVH_LAYOUT_SPLIT16),
                          &cb));
    TDVH_TRY(launch_bt_split(TL, TR, width, height, depth, (uint32_t *)cb, s));
    tdvh_dsi cd;
    TDVH_TRY(tdvh_dsi_init(&cd, cb, width, height, depth, TDVH_U16, TDVH_LAYOUT_SPLIT16));
    return tdvh_sgm(ctx, &cd, params, nullptr, disp, disp_pitch, stream);
}
